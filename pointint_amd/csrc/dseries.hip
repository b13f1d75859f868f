// Affine derivative columns dA/dtheta on the moment-series shells of triangle
// objects (AffineDoF), for the Gauss-Newton Hessian and the gradients of the
// one-pass derivative path: the derivative of the affine triangle form factor
// (cpp/src/dof/affine_dof.cpp:171-232) taken through the moment series of
// moments.hip instead of face by face.
//
// With p = A^T d, m = cof(A)^T d (cof(A) = det(A) A^-T maps reference area
// vectors to world ones), reference corners u_fj = vref_fj - c_ref and
// reference area vectors S_f, the forward series reads
//
//   A(kappa d) = e^{i kappa d.c} sum_n (i kappa)^n 2/(n+2)! G_n,   c = A c_ref + t,
//   G_n        = sum_f (m.S_f) h_n(p.u_f0, p.u_f1, p.u_f2)
//
// (p.u_fj = d.(v_fj - c) and m.S_f = d.S'_f in world terms).  Its derivatives:
//
//   dA/dt_r  = i kappa d_r A,
//   dA/dA_rc = e^{i kappa d.c} [ i kappa d_r c_ref,c S_G + d_r S_T,c
//                                + S_M,c-1 (a_c+1 x d)_r + S_M,c+1 (d x a_c+2)_r ],
//
// with a_c the columns of A (indices mod 3), S_X = sum_n (i kappa)^n 2/(n+2)! X_n
// and the vector sums
//
//   T_n = sum_f (m.S_f) sum_j dh_n/dw_j u_fj     (dG_n/dp),
//   M_n = sum_f h_n(w_f) S_f                     (dG_n/dm),
//
// where dh_n/dw_j = h_{n-1}(w_f0, w_f1, w_f2, w_fj), the complete homogeneous
// polynomial with w_j repeated.  Euler's relation for the degree-n h_n gives
// p.T_n = n G_n, and m.M_n = G_n, so only the components normal to p and m
// need sums of their own:
//
//   T_n = (n G_n / |p|^2) p + T1_n e1 + T2_n e2    (e1, e2 unit, normal to p),
//   M_n = (G_n / |m|^2) m + M1_n f1 + M2_n f2      (f1, f2 unit, normal to m).
//
// Five families of per-direction sums, one FP64 pass over the faces each:
// G_n and M1_n, M2_n (4 operations per order and face: the h_n recurrence of
// moments.hip and one FMA), T1_n, T2_n (9: h_n, the three h_{n-1}(w, w_j) by
// h_m(w, w_j) = w_j h_{m-1}(w, w_j) + h_m(w), three FMAs).  k_dseries_tri
// writes their raw sums per (family, face split, object, 64 directions);
// k_dseries_eval sums the splits and evaluates each family's series at every
// series shell (family 0 also sum_n n (i kappa)^n 2/(n+2)! G_n, for T);
// k_dseries_columns forms the 12 columns.  Order and reach are the object's
// forward series ones; the derivative terms carry an extra factor n, so their
// tail bound is (P+1) times the forward one.
#include "device_common.hpp"
#include "dtheta_columns.hpp"
#include "pii_internal.hpp"
#include "series_common.hpp"

namespace pii {

namespace {

using ser::for_chunked;
using ser::kLanes;
using ser::kSeries;
using ser::kWaves;
using ser::P;
using ser::store_split;
using ser::sum_waves_123;

constexpr int kFamilies = 5;   // G, M1, M2, T1, T2
constexpr int kSlots = 6;      // their series, then sum n (i kappa)^n c_n G_n
constexpr int kRecDser = 12;   // per face: u_0, u_1, u_2 (vref - c_ref), S (reference)

// Per face of a derivative-series object: reference corners about the
// reference centre and the reference area vector.
__global__ void __launch_bounds__(256) k_dseries_records(const ObjDesc* __restrict__ objs,
                                                         const int* __restrict__ eprefix, int N,
                                                         int total, double* __restrict__ out,
                                                         const long long* __restrict__ ooff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int o = find_owner(eprefix, N, i);
  const ObjDesc& d = objs[o];
  if (d.type != PII_GEOM_TRIANGLE || !d.dser) return;
  const int f = i - eprefix[o];
  const int i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
  const double* V = d.verts;
  double* r = out + ooff[o] + (size_t)kRecDser * f;
  double e1[3], e2[3];
  for (int c = 0; c < 3; ++c) {
    r[c] = V[3 * i0 + c] - d.rctr[c];
    r[3 + c] = V[3 * i1 + c] - d.rctr[c];
    r[6 + c] = V[3 * i2 + c] - d.rctr[c];
    e1[c] = V[3 * i1 + c] - V[3 * i0 + c];
    e2[c] = V[3 * i2 + c] - V[3 * i0 + c];
  }
  r[9] = 0.5 * (e1[1] * e2[2] - e1[2] * e2[1]);
  r[10] = 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]);
  r[11] = 0.5 * (e1[0] * e2[1] - e1[1] * e2[0]);
}

// Unit vectors a, b normal to v (and to each other): a = v x axis / |.|,
// axis the coordinate axis least aligned with v; b = v x a / |v|.
__device__ __forceinline__ void unit_normals(const double (&v)[3], double (&a)[3],
                                             double (&b)[3]) {
  const double ax = fabs(v[0]), ay = fabs(v[1]), az = fabs(v[2]);
  double c[3];
  if (ax <= ay && ax <= az) {         // axis x: v x e_x = (0, v_z, -v_y)
    c[0] = 0.0; c[1] = v[2]; c[2] = -v[1];
  } else if (ay <= az) {              // axis y: v x e_y = (-v_z, 0, v_x)
    c[0] = -v[2]; c[1] = 0.0; c[2] = v[0];
  } else {                            // axis z: v x e_z = (v_y, -v_x, 0)
    c[0] = v[1]; c[1] = -v[0]; c[2] = 0.0;
  }
  const double ic = 1.0 / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int k = 0; k < 3; ++k) a[k] = c[k] * ic;
  const double iv = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  b[0] = (v[1] * a[2] - v[2] * a[1]) * iv;
  b[1] = (v[2] * a[0] - v[0] * a[2]) * iv;
  b[2] = (v[0] * a[1] - v[1] * a[0]) * iv;
}

// Per-direction frame: p = A^T d, m = cof(A)^T d and the unit normals of
// both (A row-major in dofs[3:12]; column c of A is (A_0c, A_1c, A_2c)).
struct Frame {
  double p[3], m[3], e1[3], e2[3], f1[3], f2[3];
};
template <class PtrT>
__device__ __forceinline__ Frame dser_frame(PtrT A, double dx, double dy, double dz) {
  Frame F;
  for (int c = 0; c < 3; ++c) F.p[c] = A[c] * dx + A[3 + c] * dy + A[6 + c] * dz;
  // m_a = d . (a_{a+1} x a_{a+2})
  for (int a = 0; a < 3; ++a) {
    const int b = (a + 1) % 3, e = (a + 2) % 3;
    const double cx = A[3 + b] * A[6 + e] - A[6 + b] * A[3 + e];
    const double cy = A[6 + b] * A[e] - A[b] * A[6 + e];
    const double cz = A[b] * A[3 + e] - A[3 + b] * A[e];
    F.m[a] = dx * cx + dy * cy + dz * cz;
  }
  unit_normals(F.p, F.e1, F.e2);
  unit_normals(F.m, F.f1, F.f2);
  return F;
}

// One workgroup per (object, 64 directions, family, face split): the four
// waves take every fourth face of the split; the sums of the four waves are
// stored raw (orders 0..PP) at slot (((fam * nsplit + split) * N + obj) *
// nblk + blk) * (P + 1) * 64.
template <int PP, int FAM>
__device__ __forceinline__ void dseries_body(const ObjDesc& d, BlockObj bo, int split, int nsplit,
                                             double* red_lds, const double* __restrict__ krep,
                                             int Qh, int Mh, const double* __restrict__ srec,
                                             long long soff, double* __restrict__ part) {
  double (&red)[PP + 1][kLanes] = *reinterpret_cast<double (*)[PP + 1][kLanes]>(red_lds);
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = bo.blk * kLanes + lane;
  const int jc = j < Mh ? j : Mh - 1;
  const double dx = krep[3 * (size_t)Qh + jc], dy = krep[4 * (size_t)Qh + jc],
               dz = krep[5 * (size_t)Qh + jc];
  const Frame F = dser_frame(as_const(d.dofs) + 3, dx, dy, dz);
  // projection of the area vector (families 0-2) or of the corners (3, 4)
  double pv[3];
  for (int k = 0; k < 3; ++k)
    pv[k] = FAM == 0 ? F.m[k] : FAM == 1 ? F.f1[k] : FAM == 2 ? F.f2[k] : FAM == 3 ? F.e1[k]
                                                                                   : F.e2[k];
  double H[PP + 1];
#pragma unroll
  for (int n = 0; n <= PP; ++n) H[n] = 0.0;
  const auto* rec = as_const(srec + soff);
  const int chunk = (d.n + nsplit - 1) / nsplit;
  const int fb = split * chunk;
  const int fe = min(d.n, fb + chunk);
  double t[kRecDser] = {};
  if (fb < fe) {
    const auto* r = rec + (size_t)kRecDser * min(fb + wave, fe - 1);
#pragma unroll
    for (int i = 0; i < kRecDser; ++i) t[i] = r[i];
  }
  for (int f = fb + wave; f < fe; f += kWaves) {
    const double w0 = F.p[0] * t[0] + F.p[1] * t[1] + F.p[2] * t[2];
    const double w1 = F.p[0] * t[3] + F.p[1] * t[4] + F.p[2] * t[5];
    const double w2 = F.p[0] * t[6] + F.p[1] * t[7] + F.p[2] * t[8];
    double beta = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (FAM <= 2) {
      beta = pv[0] * t[9] + pv[1] * t[10] + pv[2] * t[11];
    } else {
      const double s = F.m[0] * t[9] + F.m[1] * t[10] + F.m[2] * t[11];
      g0 = s * (pv[0] * t[0] + pv[1] * t[1] + pv[2] * t[2]);
      g1 = s * (pv[0] * t[3] + pv[1] * t[4] + pv[2] * t[5]);
      g2 = s * (pv[0] * t[6] + pv[1] * t[7] + pv[2] * t[8]);
    }
    // the next face's record is loaded before this face's sums
    const auto* rn = rec + (size_t)kRecDser * min(f + kWaves, fe - 1);
#pragma unroll
    for (int i = 0; i < kRecDser; ++i) t[i] = rn[i];
    __builtin_amdgcn_sched_barrier(0);
    double b = 1.0, hab = 1.0, habc = 1.0;  // w1^n, h_n(w0, w1), h_n(w0, w1, w2)
    if (FAM <= 2) {
      H[0] += beta;
#pragma unroll
      for (int n = 1; n <= PP; ++n) {
        b *= w1;
        hab = fma(w0, hab, b);
        habc = fma(w2, habc, hab);
        H[n] = fma(beta, habc, H[n]);
      }
    } else {
      // G_j = h_{n-1}(w, w_j), from h_0(w, w_j) = 1
      double G0 = 1.0, G1 = 1.0, G2 = 1.0;
#pragma unroll
      for (int n = 1; n <= PP; ++n) {
        H[n] = fma(g0, G0, fma(g1, G1, fma(g2, G2, H[n])));
        if (n < PP) {
          b *= w1;
          hab = fma(w0, hab, b);
          habc = fma(w2, habc, hab);
          G0 = fma(w0, G0, habc);
          G1 = fma(w1, G1, habc);
          G2 = fma(w2, G2, habc);
        }
      }
    }
  }
  double* slot = part + ((((size_t)FAM * nsplit + split) * gridDim.y + bo.obj) * gridDim.x +
                         bo.blk) * (size_t)(P + 1) * kLanes + lane;
  store_split<PP>(H, red, wave, lane, slot);
}

__global__ void __launch_bounds__(kLanes * kWaves) k_dseries_tri(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    const double* __restrict__ srec, const long long* __restrict__ soff, double* __restrict__ part,
    int nsplit) {
  __shared__ double red[P + 1][kLanes];
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  if (!d.dser) return;  // uniform over the workgroup
  const int fam = (int)blockIdx.z / nsplit, split = (int)blockIdx.z % nsplit;
  const long long so = soff[bo.obj];
#define PII_DSER_FAM(PP, FAM) \
  dseries_body<PP, FAM>(d, bo, split, nsplit, &red[0][0], krep, Qh, Mh, srec, so, part)
#define PII_DSER_BODY(PP)                     \
  switch (fam) {                              \
    case 0: PII_DSER_FAM(PP, 0); break;       \
    case 1: PII_DSER_FAM(PP, 1); break;       \
    case 2: PII_DSER_FAM(PP, 2); break;       \
    case 3: PII_DSER_FAM(PP, 3); break;       \
    default: PII_DSER_FAM(PP, 4); break;      \
  }
  PII_SERIES_SWITCH(PII_DSER_BODY)
#undef PII_DSER_BODY
#undef PII_DSER_FAM
}

// One workgroup per (object, 64 directions, family): the waves sum every
// fourth split, the sums are reduced and scaled by c_n = (-1)^floor(n/2)
// 2/(n+2)! in LDS, then each wave evaluates E + i kappa O (even and odd
// parts in kappa^2) at every fourth series shell and stores it to
// sval[voff + fam * qbeg + node]; family 0 also stores the series of n c_n G_n
// to slot 5.
template <int PP>
__device__ __forceinline__ void dseries_eval_body(const ObjDesc& d, BlockObj bo, int fam,
                                                  double* red_lds, const double* __restrict__ krep,
                                                  int Qh, int Mh, const double* __restrict__ part,
                                                  int nsplit, double2* __restrict__ sval) {
  double (&red)[PP + 1][kLanes] = *reinterpret_cast<double (*)[PP + 1][kLanes]>(red_lds);
  const int lane = threadIdx.x & (kLanes - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = bo.blk * kLanes + lane;
  const int jc = j < Mh ? j : Mh - 1;
  double H[PP + 1];
#pragma unroll
  for (int n = 0; n <= PP; ++n) H[n] = 0.0;
  const size_t stride = (size_t)gridDim.y * gridDim.x * (P + 1) * kLanes;  // per split
  const double* src = part + (((size_t)fam * nsplit * gridDim.y + bo.obj) * gridDim.x + bo.blk) *
                                 (size_t)(P + 1) * kLanes;
  for (int z = wave; z < nsplit; z += kWaves) {
    const double* o = src + (size_t)z * stride;
    // all orders of a split in flight at once (a latency-bound sum)
#pragma unroll
    for (int n = 0; n <= PP; ++n) H[n] += o[n * kLanes + lane];
  }
  sum_waves_123<PP>(H, red, wave, lane);
  if (wave == 0)
    for_chunked<PP>([&](int n) { red[n][lane] = (red[n][lane] + H[n]) * kSeries.c[n]; });
  __syncthreads();
  const int nnode = d.qbeg;
  for (int s = wave; s < d.nser; s += kWaves) {
    const int q = s * Mh + jc;
    const double kap = krep[6 * (size_t)Qh + q];
    const double k2 = kap * kap;
    constexpr int ne = PP / 2, no = (PP - 1) / 2;
    double E = red[2 * ne][lane], O = red[2 * no + 1][lane];
    double En = (2.0 * ne) * E, On = (2.0 * no + 1.0) * O;
#pragma unroll
    for (int t = ne - 1; t >= 0; --t) {
      const double a = red[2 * t][lane];
      E = fma(E, k2, a);
      if (fam == 0) En = fma(En, k2, (2.0 * t) * a);
    }
#pragma unroll
    for (int t = no - 1; t >= 0; --t) {
      const double a = red[2 * t + 1][lane];
      O = fma(O, k2, a);
      if (fam == 0) On = fma(On, k2, (2.0 * t + 1.0) * a);
    }
    if (j < Mh) {
      const int node = s * Mh + j;
      sval[(size_t)fam * nnode + node] = make_double2(E, kap * O);
      if (fam == 0) sval[(size_t)kFamilies * nnode + node] = make_double2(En, kap * On);
    }
  }
}

__global__ void __launch_bounds__(kLanes * kWaves) k_dseries_eval(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    const double* __restrict__ part, int nsplit, double2* __restrict__ sval,
    const long long* __restrict__ voff) {
  __shared__ double red[P + 1][kLanes];
  const BlockObj bo = xcd_block_map();
  const ObjDesc& d = objs[bo.obj];
  if (!d.dser) return;
  const int fam = blockIdx.z;
  double2* sv = sval + voff[bo.obj];
#define PII_DEVAL_BODY(PP) \
  dseries_eval_body<PP>(d, bo, fam, &red[0][0], krep, Qh, Mh, part, nsplit, sv)
  PII_SERIES_SWITCH(PII_DEVAL_BODY)
#undef PII_DEVAL_BODY
}

// One wave per (object, 64 series nodes): the 12 columns sqrt(w) dA/dtheta_t
// of each node from its six series values.
__global__ void __launch_bounds__(kLanes) k_dseries_columns(
    const ObjDesc* __restrict__ objs, const double* __restrict__ krep, int Qh, int Mh,
    const double2* __restrict__ sval, const long long* __restrict__ voff, float2* __restrict__ Xt,
    int ldX, unsigned* __restrict__ colmax) {
  const int obj = blockIdx.y, lane = threadIdx.x;
  const ObjDesc& d = objs[obj];
  const int q0 = blockIdx.x * kLanes;
  if (!d.dser || q0 >= d.qbeg) return;  // uniform over the wave
  const int q = q0 + lane;
  const bool active = q < d.qbeg;
  const int qc = active ? q : d.qbeg - 1;
  const double kx = krep[qc], ky = krep[(size_t)Qh + qc], kz = krep[2 * (size_t)Qh + qc];
  const double dv[3] = {krep[3 * (size_t)Qh + qc], krep[4 * (size_t)Qh + qc],
                        krep[5 * (size_t)Qh + qc]};
  const double kap = krep[6 * (size_t)Qh + qc];
  const auto* A = as_const(d.dofs) + 3;
  const Frame F = dser_frame(A, dv[0], dv[1], dv[2]);
  const int nnode = d.qbeg;
  const double2* sv = sval + voff[obj] + qc;
  const double2 SG = sv[0], SM1 = sv[(size_t)nnode], SM2 = sv[2 * (size_t)nnode],
                ST1 = sv[3 * (size_t)nnode], ST2 = sv[4 * (size_t)nnode],
                SGn = sv[5 * (size_t)nnode];
  const double ip = 1.0 / (F.p[0] * F.p[0] + F.p[1] * F.p[1] + F.p[2] * F.p[2]);
  const double im = 1.0 / (F.m[0] * F.m[0] + F.m[1] * F.m[1] + F.m[2] * F.m[2]);
  double2 ST[3], SM[3];
  for (int c = 0; c < 3; ++c) {
    const double tp = F.p[c] * ip, mp = F.m[c] * im;
    ST[c] = make_double2(SGn.x * tp + ST1.x * F.e1[c] + ST2.x * F.e2[c],
                         SGn.y * tp + ST1.y * F.e1[c] + ST2.y * F.e2[c]);
    SM[c] = make_double2(SG.x * mp + SM1.x * F.f1[c] + SM2.x * F.f2[c],
                         SG.y * mp + SM1.y * F.f1[c] + SM2.y * F.f2[c]);
  }
  double sn, cs;
  sincos(kap * (dv[0] * d.ctr[0] + dv[1] * d.ctr[1] + dv[2] * d.ctr[2]), &sn, &cs);
  double v[20];
  v[0] = cs * SG.x - sn * SG.y;
  v[1] = sn * SG.x + cs * SG.y;
  for (int c = 0; c < 3; ++c) {
    // columns a_{c+1}, a_{c+2} of A
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    const double a1[3] = {A[c1], A[3 + c1], A[6 + c1]};
    const double a2[3] = {A[c2], A[3 + c2], A[6 + c2]};
    const double X1[3] = {a1[1] * dv[2] - a1[2] * dv[1], a1[2] * dv[0] - a1[0] * dv[2],
                          a1[0] * dv[1] - a1[1] * dv[0]};  // a_{c+1} x d
    const double X2[3] = {dv[1] * a2[2] - dv[2] * a2[1], dv[2] * a2[0] - dv[0] * a2[2],
                          dv[0] * a2[1] - dv[1] * a2[0]};  // d x a_{c+2}
    for (int r = 0; r < 3; ++r) {
      // i kappa d_r c_ref,c S_G + d_r S_T,c + S_M,c-1 X1_r + S_M,c+1 X2_r
      const double ph = kap * dv[r] * d.rctr[c];
      const double re = -ph * SG.y + dv[r] * ST[c].x + SM[c2].x * X1[r] + SM[c1].x * X2[r];
      const double imv = ph * SG.x + dv[r] * ST[c].y + SM[c2].y * X1[r] + SM[c1].y * X2[r];
      const int idx = 2 + 2 * (3 * r + c);
      v[idx] = cs * re - sn * imv;
      v[idx + 1] = sn * re + cs * imv;
    }
  }
  if (kap < 1e-10) {  // dA/dtheta = 0 at k = 0 (affine_dof.cpp:181)
    for (int i = 0; i < 20; ++i) v[i] = 0.0;
  }
  write_columns(v, q, Qh, kx, ky, kz, krep, d.col, Xt, ldX, colmax, lane, active);
}

}  // namespace

void launch_dseries_records(hipStream_t s, const ObjDesc* d_objs, const int* d_eprefix, int N,
                            int total_e, double* d_out, const long long* d_off) {
  if (total_e <= 0) return;
  hipLaunchKernelGGL(k_dseries_records, dim3((total_e + 255) / 256), dim3(256), 0, s, d_objs,
                     d_eprefix, N, total_e, d_out, d_off);
}

size_t dseries_record_doubles(int nfaces) { return (size_t)kRecDser * nfaces; }

int dseries_slots() {
  static std::atomic<int> cache[16];
  return resident_workgroups(k_dseries_tri, kLanes * kWaves, cache);
}

size_t dseries_part_doubles(int N, int Mh, int nsplit) {
  return (size_t)kFamilies * nsplit * N * ((Mh + kLanes - 1) / kLanes) * (P + 1) * kLanes;
}

size_t dseries_value_count(int qbeg) { return (size_t)kSlots * qbeg; }

void launch_dseries(hipStream_t s, const ObjDesc* d_objs, int N, const double* d_krep, int Qh,
                    int Mh, const double* d_srec, const long long* d_soff, double* d_part,
                    int nsplit, double2* d_sval, const long long* d_voff, int max_qbeg,
                    float2* d_Xt, int ldX, unsigned* d_colmax) {
  if (N <= 0 || Mh <= 0 || max_qbeg <= 0) return;
  const int nblk = (Mh + kLanes - 1) / kLanes;
  hipLaunchKernelGGL(k_dseries_tri, dim3(nblk, N, kFamilies * nsplit), dim3(kLanes * kWaves), 0,
                     s, d_objs, d_krep, Qh, Mh, d_srec, d_soff, d_part, nsplit);
  hipLaunchKernelGGL(k_dseries_eval, dim3(nblk, N, kFamilies), dim3(kLanes * kWaves), 0, s,
                     d_objs, d_krep, Qh, Mh, d_part, nsplit, d_sval, d_voff);
  hipLaunchKernelGGL(k_dseries_columns, dim3((max_qbeg + kLanes - 1) / kLanes, N), dim3(kLanes),
                     0, s, d_objs, d_krep, Qh, Mh, d_sval, d_voff, d_Xt, ldX, d_colmax);
}

}  // namespace pii
